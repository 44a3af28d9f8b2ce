"""DirectionalLight on the device against the oracle (which is held to the reference's own
DirectionalLight and PathIntegrator in tests/test_oracle_directional.py): hard and cone-jittered suns
beside a dome or an area light, radiance within 1e-3 relative on every pixel at identical sample
indices, image means within 1e-3; replay / adjoint refuse the delta light."""
import numpy as np
import pytest

from anacapa_amd import _abi as A
from test_gpu_parity import _radiance_close
from test_oracle_directional import CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(CASES))
def test_directional_light_vs_oracle(gpu, oracle, name):
    sc = CASES[name]()
    print(f"[directional] {name}", flush=True)
    gpu.prepare(sc)
    got = gpu.render_tile(0, 0, 64, 64, 0, 4)
    ref = oracle.scene(sc).render_tile(0, 0, 64, 64, 0, 4)
    _radiance_close(got, ref, what=f"directional {name}")


def test_directional_light_refused_by_replay(gpu):
    from anacapa_amd.integrator import BackendError
    gpu.prepare(CASES["cornell_cone1"]())
    with pytest.raises(BackendError) as e:
        gpu.replay(0, 1)
    assert e.value.status == A.AHIP_E_UNSUPPORTED
