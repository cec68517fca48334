"""The autotuner: picks a setting among its candidates, caches it per size,
and the engine keeps oracle parity under the setting it picked."""
import numpy as np
import pytest

import oracle as O
from helpers import to_oracle

pytestmark = pytest.mark.gpu


def test_autotune_setting_cache_and_parity(capi, tmp_path, monkeypatch):
    from pennylane_lightning_kokkos_amd import autotune

    for k in autotune.KNOBS:
        monkeypatch.delenv(k, raising=False)
    cache = str(tmp_path / "tune.json")
    best = autotune.tune(19, np.complex128, n_obs=3, layers=2, repeats=1, cache=cache)
    if best["PLK_RTILE"] == "1":
        assert best["PLK_RTILE_NJ"] in {"2", "3", "4"}
    else:
        assert best["PLK_RTILE"] == "0"
        assert best["PLK_TILE_BITS_ADJ"] in {"9", "10", "11"}
        assert best["PLK_TILE_NJ"] in {"2", "3", "4"}
        assert best["PLK_TILE_BITS_FWD"] in {"9", "10", "11"}
    assert autotune.tune(19, np.complex128, n_obs=3, cache=cache) == best  # cache hit
    for k, v in best.items():
        monkeypatch.setenv(k, v)

    names, params, wires = autotune._circuit(19, 2, seed=19)
    inv, mats = [False] * len(names), [[] for _ in names]
    obs = [("named", "PauliZ", [w]) for w in range(3)]
    tp = list(range(sum(1 for p in params if p)))
    s = capi.State(19)
    got = capi.adjoint_jacobian(s, [capi.Obs.named("PauliZ", [w]) for w in range(3)],
                                capi.Ops(names, params, wires, inv, mats), tp, True)
    ref = O.adjoint_jacobian(O.zero_state(19), [to_oracle(o) for o in obs],
                             O.Ops(names, params, wires, inv, mats), tp, True)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)
